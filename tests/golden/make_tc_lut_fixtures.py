"""Extract the TC-LUT build cases of the reference's tests as a compact fixture.

Run once in the build container (it reads /root/reference); the output
tests/golden/tc_lut_build.json is committed, and nothing at run time reads the
reference. Every value is the reference's exact f32 word (uint32) or integer;
nothing here evaluates the algorithm under test.

Source: tests/ffi/fixtures/tc_lut/records-0.json, "leaves": the 23 bounded
build cases that tests/ffi/tc_lut_test.cpp:81-158 drives through the native
builder. Cases 0-8 are built tables (their complete 147456 f32 words are pinned
by the FNV-1a digest of their little-endian bytes, as the reference's test
pins them); cases 9-22 are failures with the builder's diagnostic. The
spectra of every case but 7 and 8 (the bundled Hanatos / Arctic resources,
not extracted) are the test's procedural table:
word(i) = 0x3e000123 + (i mod 4093) * 173 over the 192 x 192 x 81
(C, M, wavelength) samples, with sample 0 set to +inf for case 21 and to a
quiet NaN for case 22 (tc_lut_test.cpp:85-104).
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/ffi/fixtures/tc_lut/records-0.json"


def main():
    with open(SRC) as f:
        d = json.load(f)
    leaves = d["leaves"]
    hull = leaves["hull"]
    cases = []
    for row in leaves["cases"]:
        i = row["input"]
        cases.append({
            "id": i["id"], "source": i["source"], "method": i["method"], "blur": i["blur"],
            "surface": bool(i["surface"]), "compression": bool(i["compression"]),
            "hull_available": bool(i["hull_available"]),
            "sensitivity": i["sensitivity"], "spd": i["spd"], "white": i["white"],
            "surface_params": i["surface_params"],
            "built": bool(row["built"]), "diagnostic": row["diagnostic"],
            "digest": row["table"]["native_byte_digest_fnv1a64"] if row["built"] else None,
            "selected": row["table"]["selected"] if row["built"] else None,
        })
    out = {"source": "tests/ffi/fixtures/tc_lut/records-0.json (leaves)",
           "procedural": {"base_word": 0x3e000123, "period": 4093, "step": 173},
           "hull": {"center": hull["center"], "xy": hull["xy"]},
           "cases": cases}
    path = os.path.join(HERE, "tc_lut_build.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {len(cases)} cases to {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()

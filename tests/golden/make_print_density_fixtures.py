"""Extract the reference's density-model and print-gamma characterizations as
one compact fixture (tests/golden/print_density.json).

Run once in the build container (it reads /root/reference); the output is
committed, and nothing at run time reads the reference. Every value is a word
the reference's files hold (f64 or f32 bits, u64 identities); nothing here
evaluates the algorithms under test.

Sources (tests/profile/fixtures/):
  * density.json + density-samples.bin: 138 density-model cases (the 28 shipped
    profiles' models, the eight print models gamma-adjusted, synthetic edge
    models) over 11 authored f64 exposure axes, with the approved Rust words of
    every sample (libm 0.2.15 erfcf) and, for 119 complete cases, the total,
    layer and raw print-curve identities. The 19,247 sample records are kept
    here only where no identity covers them (the three cases with historical
    failure rows, which have none) plus every 64th record as direct words; the
    identities pin all the others. Records without a valid sample are listed
    per case.
  * the Linux native witnesses of density.json (Linux debug and release
    agree): the first sample where the reference's native build (glibc
    erfcf) differs from the approved words;
  * print-gamma.json: 78 captured gamma cases (the unadjusted source model; the
    adjusted model is the density case of the same id) and 17 analytical cases
    (model, axis, gamma -> totals and print identity).
"""
from __future__ import annotations

import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests/profile/fixtures"
RECORD = 49
STRIDE = 64


def read_json(name):
    with open(os.path.join(REF, name)) as f:
        return json.load(f)


def record(blob, index):
    row = blob[index * RECORD:(index + 1) * RECORD]
    valid = row[0]
    return valid, list(struct.unpack("<12I", row[1:]))


def main():
    density = read_json("density.json")
    gamma = read_json("print-gamma.json")
    with open(os.path.join(REF, "density-samples.bin"), "rb") as f:
        blob = f.read()
    assert len(blob) == 19247 * RECORD
    cases = []
    for case in density["cases"]:
        out = {k: case[k] for k in ("id", "positive", "gamma_bits", "adjusted", "axis", "supported", "model",
                                    "identities")}
        if case["supported"]:
            n = len(density["axes"][case["axis"]])
            start = case["sample_start"]
            rows = [record(blob, start + i) for i in range(n)]
            out["invalid"] = [i for i, (v, _) in enumerate(rows) if not v]
            keep_all = case["identities"] is None
            out["words"] = {str(i): w for i, (v, w) in enumerate(rows) if v and (keep_all or i % STRIDE == 0)}
        cases.append(out)
    witnesses = []
    for w in density["native_witnesses"]:
        if w["platform"] != "linux":
            continue
        case = next(c for c in density["cases"] if c["id"] == w["case"])
        valid, words = record(blob, case["sample_start"] + w["sample"])
        assert valid
        witnesses.append({"case": w["case"], "sample": w["sample"], "native": w["native"], "rust": words})
    out = {
        "axes": density["axes"],
        "cases": cases,
        "native_witnesses": witnesses,
        "gamma_captured": [{"id": c["id"], "source_model": c["source_model"]} for c in gamma["captured"]],
        "gamma_analytical": gamma["analytical"],
        "sources": ["tests/profile/fixtures/density.json", "tests/profile/fixtures/density-samples.bin",
                    "tests/profile/fixtures/print-gamma.json"],
    }
    path = os.path.join(HERE, "print_density.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes): {len(cases)} density cases, "
          f"{len(witnesses)} native witnesses, {len(out['gamma_captured'])} + {len(out['gamma_analytical'])} gamma")


if __name__ == "__main__":
    main()

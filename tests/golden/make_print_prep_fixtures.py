"""Extract the reference's print-preparation characterizations as one compact
fixture (tests/golden/print_prep.json).

Run once in the build container (it reads /root/reference); the output is
committed, and nothing at run time reads the reference. Every value is the
reference's exact f32 word (uint32) or integer; nothing here evaluates the
algorithms under test. Word arrays of 81 or more entries are stored once in
"arrays" and referenced from the cases as {"@": index}.

Sources (linux-release presets, the native expectations):
  * tests/print/illuminant-fixtures/linux-release-rust-literal.json: 127
    filtered print illuminant cases (dichroic edges and transitions, CMY cc,
    source SPD -> the 81 filtered samples, or the failure diagnostic), the
    words of the reference's Rust preparation with its pinned libm 0.2.15
    erff (the file print.rs:798 tests against; the native C++ capture beside
    it differs in two words, from glibc's erff);
  * tests/print/preflash-fixtures/linux-release.json: 19 preflash integral
    cases (film base density, filtered illuminant, print sensitivity -> the
    raw CMY words, or the failure diagnostic) and, from the manifest, the
    first failing channel of each failure case;
  * tests/print/balance-fixtures/linux-release.json: 42 print balance cases
    (base and channel densities, filtered illuminant, print sensitivity,
    baseline / compensated densities, normalization mode, reference hash ->
    raw RGB, the three balance scalars and the identity hashes, or the
    failure diagnostic) and, from the manifest, each case's status category,
    failure reason and channel of the checked ABI.
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests/print"


def read(path):
    with open(os.path.join(REF, path)) as f:
        return json.load(f)


def load(path):
    return read(path)["cases"]


def pooled(sections):
    arrays, index = [], {}
    for cases in sections:
        for case in cases:
            for key, value in case.items():
                if isinstance(value, list) and len(value) >= 81:
                    slot = index.setdefault(tuple(value), len(arrays))
                    if slot == len(arrays):
                        arrays.append(value)
                    case[key] = {"@": slot}
    return arrays


def main():
    out = {
        "illuminant": load("illuminant-fixtures/linux-release-rust-literal.json"),
        "preflash": load("preflash-fixtures/linux-release.json"),
        "balance": load("balance-fixtures/linux-release.json"),
        "preflash_failure_channels": read("preflash-fixtures/manifest.json")["failure_channels"],
        "balance_outcomes": read("balance-fixtures/manifest.json")["outcomes"],
        "sources": ["tests/print/illuminant-fixtures/linux-release-rust-literal.json",
                    "tests/print/preflash-fixtures/linux-release.json",
                    "tests/print/balance-fixtures/linux-release.json",
                    "tests/print/preflash-fixtures/manifest.json",
                    "tests/print/balance-fixtures/manifest.json"],
    }
    out["arrays"] = pooled([out["illuminant"], out["preflash"], out["balance"]])
    path = os.path.join(HERE, "print_prep.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes): "
          f"{len(out['illuminant'])} illuminant, {len(out['preflash'])} preflash, {len(out['balance'])} balance cases")


if __name__ == "__main__":
    main()

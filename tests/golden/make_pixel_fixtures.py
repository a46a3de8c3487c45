"""Extract compact per-pixel golden fixtures from the reference's test data.

Run once in the build container (it reads /root/reference); the outputs under
tests/golden/pixel/ are committed, and nothing at run time reads the
reference. Every array holds the reference's exact f32 words (uint32) or its
exact integers; nothing here evaluates the algorithms under test.

Sources:
  * tests/ffi/fixtures/exposure/mallett-numerical-linux-release.json
      -> mallett_focused.npz: the 96 "focused" Mallett mid-gray cases
      (input matrices, flags, basis/illuminant/sensitivity; expected working
      XYZ, linear sRGB, raw B/G/R and the green normalization) computed at the
      fixed input RGB 0.184 (rust/film-juicer-core/src/exposure.rs:391-406).
  * tests/ffi/fixtures/tc_lut/records-0.json + records-0-tables.bin
      -> tc_sampling.npz: the 15 host-exposure samples of each of the 9 built
      TC LUTs (XYZ in, brightness x Mitchell-sampled RGB out;
      reconstruction.rs:684-703) with the table cells the samples touch, and
      the table's "selected" cells as a cross-check of the table words.
  * tests/print/balance-fixtures/linux-release.json
      -> print_balance.npz: the 42 literal print-balance cases (film base and
      channel densities, filtered illuminant, print sensitivity; baseline and
      compensated densities -> raw RGB and the three balance scalars;
      print.rs:435-522).
  * tests/optics/gaussian-fixtures/linux-payload.json
      -> gaussian_cases.npz: 188 Gaussian FIR weight arrays (radius, sigma
      bits, exact weights; optics/gaussian.rs:20-56).
  * tests/scanner/output_encoding_native.inc
      -> encoding_entries.npz: the nine output-space catalogue entries (CCTF
      kind, gamma/a/b/c/d/linear-cutoff words, DWG-to-output matrix words).
  * tests/ffi/fixtures/output_gamut/{records-0.json,tables-0.bin}
      -> gamut_tables.npz: three output spaces' 64x720 chroma-boundary tables
      with their native<->D65 XYZ matrices.
"""
from __future__ import annotations

import json
import os

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "pixel")


def u32(v) -> np.ndarray:
    return np.asarray(v, dtype=np.uint64).astype(np.uint32)


def mallett_focused() -> None:
    d = json.load(open(f"{REF}/tests/ffi/fixtures/exposure/mallett-numerical-linux-release.json"))
    rows = d["focused"]
    out = {
        "id": np.array([r["id"] for r in rows], np.int64),
        "space": np.array([r["input"]["space"] for r in rows], np.int32),
        "decode": np.array([r["input"]["decode"] for r in rows], np.int32),
        "adapt": np.array([r["input"]["adapt"] for r in rows], np.int32),
    }
    for key in ("rgb_to_xyz", "xyz_adapt", "xyz_to_srgb", "basis", "illuminant", "sensitivity"):
        out[key] = np.stack([u32(r["input"][key]) for r in rows])
    for key in ("linear_xyz", "linear_srgb", "raw_bgr", "normalization", "pre_xyz"):
        out["expected_" + key] = np.stack([u32(r["expected"][key]) for r in rows])
    np.savez_compressed(os.path.join(OUT, "mallett_focused.npz"), **out)


def tc_sampling() -> None:
    d = json.load(open(f"{REF}/tests/ffi/fixtures/tc_lut/records-0.json"))
    blob = np.fromfile(f"{REF}/tests/ffi/fixtures/tc_lut/records-0-tables.bin", dtype="<u4")
    cases, cells, words, xyz, expected, tc, selected = [], [], [], [], [], [], []
    for case in d["leaves"]["cases"]:
        table = case.get("table")
        if not case.get("built") or not table or "byte_offset" not in table or "samples" not in case:
            continue
        start = table["byte_offset"] // 4
        t = blob[start:start + table["count"]].reshape(192 * 192, 4)
        cid = int(case["input"]["id"])
        for sel in table["selected"]:
            selected.append([cid, sel["cell"], *[int(w) for w in sel["rgb"]]])
        need = set()
        for s in case["samples"]:
            if s.get("classification") != "native_defined":
                continue
            cases.append(cid)
            xyz.append(u32(s["xyz"]))
            expected.append(u32(s["expected"]))
            tc.append(u32(s["tc"]))
            # Every cell within 3 of the sample's square (a superset of the
            # reflected 4x4 Mitchell footprint).
            q = np.asarray(tc[-1], dtype=np.uint32).view(np.float32).astype(np.float64)
            b = np.clip(np.floor(np.clip(q, 0, 1) * 191).astype(int), 0, 191)
            for c in range(max(0, b[0] - 3), min(192, b[0] + 4)):
                for m in range(max(0, b[1] - 3), min(192, b[1] + 4)):
                    need.add(c * 192 + m)
        for cell in sorted(need):
            cells.append([cid, cell])
            words.append(t[cell])
    np.savez_compressed(os.path.join(OUT, "tc_sampling.npz"), case=np.array(cases, np.int32),
                        xyz=np.stack(xyz), expected=np.stack(expected), tc=np.stack(tc),
                        cells=np.array(cells, np.int32), cell_words=np.stack(words),
                        selected=np.array(selected, np.uint64))


def print_balance() -> None:
    d = json.load(open(f"{REF}/tests/print/balance-fixtures/linux-release.json"))
    rows = [c for c in d["cases"] if c["success"]]
    out = {"id": np.array([r["id"] for r in rows]), "mode": np.array([r["normalization_mode"] for r in rows], np.int32)}
    for key in ("film_base_density", "film_channel_density_cmy", "filtered_main_illuminant", "print_sensitivity_rgb",
                "baseline_density_cmy", "compensated_density_cmy"):
        out[key] = np.stack([u32(r[key]) for r in rows])
    for key in ("baseline_raw_rgb", "compensated_raw_rgb", "scalars"):
        out["expected_" + key] = np.stack([u32(r["balance"][key]) for r in rows])
    np.savez_compressed(os.path.join(OUT, "print_balance.npz"), **out)


def gaussian_cases() -> None:
    d = json.load(open(f"{REF}/tests/optics/gaussian-fixtures/linux-payload.json"))
    radius, sigma, offset, words, names = [], [], [0], [], []
    for c in d["cases"]:
        radius.append(c["radius"])
        sigma.append(c["sigma_bits"])
        w = [int(x) for x in c["weights"]]
        assert len(w) == c["count"]
        words.extend(w)
        offset.append(len(words))
        names.append(c["name"])
    np.savez_compressed(os.path.join(OUT, "gaussian_cases.npz"), radius=np.array(radius, np.int64),
                        sigma_bits=u32(sigma), offset=np.array(offset, np.int64), weights=u32(words),
                        name=np.array(names))


def gamut_tables(rows=(0, 11, 14)) -> None:
    d = json.load(open(f"{REF}/tests/ffi/fixtures/output_gamut/records-0.json"))
    blob = np.fromfile(f"{REF}/tests/ffi/fixtures/output_gamut/tables-0.bin", dtype="<u4")
    out = {"tag": [], "entry": [], "exit": [], "cmax": []}
    for i in rows:
        r = d["rows"][i]
        assert r["transform"]["valid"] and r["boundary_ok"] and r["table_count"] == 64 * 720
        out["tag"].append(r["tag"])
        out["entry"].append(u32(r["transform"]["entry"]))
        out["exit"].append(u32(r["transform"]["exit"]))
        out["cmax"].append(blob[r["table_offset"] // 4:r["table_offset"] // 4 + r["table_count"]])
    np.savez_compressed(os.path.join(OUT, "gamut_tables.npz"), **{k: np.stack(v) for k, v in out.items()})


def encoding_entries() -> None:
    import re
    text = open(f"{REF}/tests/scanner/output_encoding_native.inc").read()
    kinds, params, matrices = [], [], []
    for m in re.finditer(r"\{\{static_cast<CctfKind>\((\d+)\),(.*?)\}, \{(.*?)\}\}", text):
        words = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]+)u", m.group(2))]
        mat = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]+)u", m.group(3))]
        assert len(words) == 6 and len(mat) == 9
        kinds.append(int(m.group(1)))
        params.append(words)
        matrices.append(mat)
    assert len(kinds) == 9
    np.savez_compressed(os.path.join(OUT, "encoding_entries.npz"), kind=np.array(kinds, np.int32),
                        params=u32(params), dwg_to_output=u32(matrices))


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    mallett_focused()
    tc_sampling()
    print_balance()
    gaussian_cases()
    gamut_tables()
    encoding_entries()

"""Extract compact golden fixtures from the reference's own test data.

Run once in the build container (it reads /root/reference); the outputs are
committed under tests/golden/ and nothing at run time reads the reference.

Sources (all data files held by the reference's tests):
  * tests/ffi/cuda_abi_facts.inc                       -> abi_facts.json
      x86-64 sizes/offsets/values of every FjPreparedHostData record.
  * tests/ffi/prepared_call_abi_facts.inc              -> call_abi_facts.json
      x86-64 sizes/offsets/values of the call-scope and frame-value records.
  * tests/ffi/prepared-values.c                        -> boundary_abi_facts.json
      C11 layout assertions of the diffusion-components and DIR-boundary
      import records (juicer_diffusion_api.h, juicer_dir_boundary_api.h).
  * tests/ffi/fixtures/prepared-state-projection/linux-release.json
      -> prepared/<case>.npz + prepared/<case>.json
      Complete prepared products (exact f32 words) for named recipes:
      TC LUT, film density curves, DIR axes, print spectra/curves, scanner
      spectra, output colour and gamut cmax table.
  * tests/scanner/lut-fixtures/native.json             -> scanner_lut_cases.npz
      Exact scanner PCHIP LUT words for small synthetic inputs and for the
      four production LUTs (one per route).
  * tests/scatter_halation/frame-fixtures/native-payload.json
      -> halation_frame_cases.json: resolved halation inputs with the exact
      descriptor words (kinds, radii, FIR and YVV words) the Rust frame
      builder produced, plus its f64 exp primitives.
  * tests/scatter_halation/fixtures/{manifest.json,scatter_halation_reference.bin}
      -> halation_cohort.json + halation_cohort.npz: the external operator
      cohort (source carrier planes and expected planes).
  * tests/diffusion/fixtures/diffusion_reference_v1.{json,npz}
      -> diffusion_cohort.json + diffusion_cohort.npz: the external diffusion
      operator cohort (17x19 and radius-cap images, camera and enlarger
      stages, four filter families, HDR inputs) and its family, strength and
      warmth expansion tables.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import struct
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def f32(word: int) -> float:
    return struct.unpack("<f", struct.pack("<I", word & 0xFFFFFFFF))[0]


def abi_facts(source: str = "cuda_abi_facts.inc", target: str = "abi_facts.json") -> None:
    facts = {"types": {}, "fields": [], "values": {}}
    pat = re.compile(r"^FJ_ABI_(TYPE|FIELD|VALUE)\((.*)\)\s*$")
    for line in open(f"{REF}/tests/ffi/{source}"):
        m = pat.match(line.strip())
        if not m:
            continue
        kind, args = m.group(1), [a.strip() for a in m.group(2).split(",")]
        if kind == "TYPE":
            facts["types"][args[0]] = [int(args[1]), int(args[2])]
        elif kind == "FIELD":
            facts["fields"].append([args[0], args[1], int(args[2])])
        else:
            facts["values"][args[0]] = int(args[1], 0)
    with open(os.path.join(HERE, target), "w") as fh:
        json.dump(facts, fh, indent=0, sort_keys=True)


# ---------------------------------------------------------------------------
# prepared-state projection
# ---------------------------------------------------------------------------
CASES = [
    "NegativeDirectHanatos",
    "NegativePrintHanatos",
    "PositivePrintHanatos",
    "CameraAndEnlargerOptics",
    "GammaAndGrainLayers",
    "DirAndCompressionOff",
    "PreflashAndNormalizationModes[norm=1,compensate=1]",
]
# File-system friendly names for case ids that carry brackets.
ALIASES = {"PreflashAndNormalizationModes[norm=1,compensate=1]": "PrintPreflash"}


def scalar(node):
    t = node["type"]
    if t == "f32":
        return f32(int(node["bits"], 16))
    if t == "f64":
        return struct.unpack("<d", struct.pack("<Q", int(node["bits"], 16)))[0]
    if t == "utf8":
        return node["value"]
    return int(node["value"])


def array(node):
    et = node["element_type"]
    words = node["words"]
    if et == "f32":
        return np.array([int(w, 16) for w in words], dtype=np.uint32).view(np.float32)
    if et == "f64":
        return np.array([int(w, 16) for w in words], dtype=np.uint64).view(np.float64)
    if et == "u32":
        return np.array([w if isinstance(w, int) else int(w, 16) for w in words], dtype=np.uint32)
    raise ValueError(et)


def flatten(node, prefix, scalars, arrays):
    if isinstance(node, dict):
        if "words" in node and "element_type" in node:
            arrays[prefix] = array(node)
            return
        if "elements" in node and node.get("element_type", "").startswith("span"):
            for i, el in enumerate(node["elements"]):
                arrays[f"{prefix}[{i}]"] = array(el)
            return
        if "type" in node and ("value" in node or "bits" in node):
            if node["type"] in ("live-backing",):
                return
            scalars[prefix] = scalar(node)
            return
        for k, v in node.items():
            flatten(v, f"{prefix}.{k}" if prefix else k, scalars, arrays)


def prepared_cases() -> None:
    out_dir = os.path.join(HERE, "prepared")
    os.makedirs(out_dir, exist_ok=True)
    doc = json.load(open(f"{REF}/tests/ffi/fixtures/prepared-state-projection/linux-release.json"))
    shared = {}  # content-addressed large arrays, stored once
    for case in doc["cases"]:
        cid = case["case_id"]
        if cid not in CASES:
            continue
        name = ALIASES.get(cid, cid)
        scalars, arrays = {}, {}
        flatten(case["expected"], "", scalars, arrays)
        inputs = {}
        flatten(case["input"], "", inputs, {})
        refs = {}
        small = {}
        for k, v in arrays.items():
            if v.size >= 4096:
                key = f"t{v.size}_{hashlib.sha1(v.tobytes()).hexdigest()[:12]}"
                shared[key] = v
                refs[k] = key
            else:
                small[k] = v
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), **small)
        with open(os.path.join(out_dir, f"{name}.json"), "w") as fh:
            json.dump({"case_id": cid, "scalars": scalars, "large": refs, "input": inputs},
                      fh, indent=0, sort_keys=True)
    np.savez_compressed(os.path.join(out_dir, "shared_tables.npz"), **shared)


def scanner_luts() -> None:
    doc = json.load(open(f"{REF}/tests/scanner/lut-fixtures/native.json"))
    out = {}
    meta = []
    for group in ("cases", "products"):
        for case in doc[group]:
            cid = case["id"]
            inp = case["input"]
            nat = case["native"]
            entry = {"id": cid, "group": group, "success": bool(nat["success"]),
                     "resolution": int(inp["resolution"]), "medium": int(inp["medium"]),
                     "sample_count": int(inp["sample_count"]), "has_baseline": bool(inp["has_baseline"]),
                     "inv_yn": int(inp["inv_yn"])}
            if group == "products":
                entry["descriptor_hash"] = int(case["descriptor_hash"])
                entry["route"] = int(case["route"])
            meta.append(entry)
            for k in ("eps_c", "eps_m", "eps_y", "weighted_x", "weighted_y", "weighted_z",
                      "baseline", "range_min_cmy", "inverse_span_cmy"):
                out[f"{cid}/in/{k}"] = np.array(inp[k], dtype=np.uint32)
            if nat["success"]:
                for k in ("log2_xyz", "slope_c", "slope_m", "slope_y", "cell_min", "cell_max"):
                    out[f"{cid}/out/{k}"] = np.array(nat[k], dtype=np.uint32)
    slopes = []
    for s in doc["slopes"]:
        slopes.append({"id": s["id"], "input": [str(w) for w in s["input"]],
                       "native": [str(w) for w in s["native"]]})
    np.savez_compressed(os.path.join(HERE, "scanner_lut_cases.npz"), **out)
    with open(os.path.join(HERE, "scanner_lut_cases.json"), "w") as fh:
        json.dump({"cases": meta, "slopes": slopes}, fh, indent=0)


def dir_fixtures() -> None:
    """Spatial-DIR frame math and reflected-YVV boundary words
    (tests/dir/frame-fixtures, tests/dir/boundary-fixtures of the reference)."""
    frame = json.load(open(f"{REF}/tests/dir/frame-fixtures/linux-release.json"))
    rows = []
    for group in ("cases", "products"):
        for c in frame[group]:
            rows.append({"group": group, "input": c["input"], "native": c["native"]})
    with open(os.path.join(HERE, "dir_frame_cases.json"), "w") as fh:
        json.dump(rows, fh)
    bound = json.load(open(f"{REF}/tests/dir/boundary-fixtures/linux-release.json"))
    keep = []
    for c in bound["cases"]:
        if not c["success"]:
            continue
        keep.append({"input": c["input"], "horizontal": c["horizontal"], "vertical": c["vertical"]})
    with open(os.path.join(HERE, "dir_boundary_cases.json"), "w") as fh:
        json.dump(keep, fh)


def _gauss_words(g):
    return [g["kind"], g["radius"]] + list(g["weights"]) + list(g["yvv"])


def halation_fixtures() -> None:
    """Halation frame descriptors and the operator cohort."""
    base = f"{REF}/tests/scatter_halation"
    native = json.load(open(f"{base}/frame-fixtures/native-payload.json"))
    rows = []
    for group in ("cases", "products"):
        for c in native[group]:
            d = c["descriptor"]
            row = {"group": group, "input": c["input"], "success": c["success"], "present": c["present"]}
            if d is not None:
                row["descriptor"] = {
                    "recipe_hash": d["recipe_hash"], "scatter_amount": d["scatter_amount"],
                    "channels": [{"strength": ch["strength"], "core": _gauss_words(ch["core"]),
                                  "tail": [_gauss_words(g) for g in ch["tail"]],
                                  "bounce": [_gauss_words(g) for g in ch["bounce"]]} for ch in d["channels"]]}
            rows.append(row)
    with open(os.path.join(HERE, "halation_frame_cases.json"), "w") as fh:
        json.dump({"rows": rows, "primitives": native["primitives"]}, fh)

    man = json.load(open(f"{base}/fixtures/manifest.json"))
    blob = open(f"{base}/fixtures/scatter_halation_reference.bin", "rb").read()
    arrays, meta = {}, []
    for r in man["rows"]:
        entry = {"id": r["id"], "category": r["category"], "comparison": r["comparison"],
                 "recipe": r["recipeInput"]}
        for key in ("source", "expected", "downstreamExpected"):
            if key not in r:
                continue
            planes = []
            for pl in r[key]["planes"]:
                dt = np.dtype(pl["dtype"])
                a = np.frombuffer(blob, dtype=dt, count=pl["elementCount"], offset=pl["byteOffset"])
                assert hashlib.sha256(a.tobytes()).hexdigest() == pl["sha256"], (r["id"], key)
                planes.append(a.reshape(pl["shape"]))
            arrays[f"{r['id']}/{key}"] = np.stack(planes)
            entry.setdefault("planes", []).append(key)
        meta.append(entry)
    np.savez_compressed(os.path.join(HERE, "halation_cohort.npz"), **arrays)
    with open(os.path.join(HERE, "halation_cohort.json"), "w") as fh:
        json.dump(meta, fh, indent=0)


def diffusion_fixtures() -> None:
    """Diffusion operator cohort (inputs/outputs only; PSF arrays dropped)."""
    base = f"{REF}/tests/diffusion/fixtures/diffusion_reference_v1"
    man = json.load(open(base + ".json"))
    z = np.load(base + ".npz", allow_pickle=False)
    keep = {}
    meta = {k: man[k] for k in ("family_order", "family_cases", "strength_cases", "warmth_cases")}
    for group in ("image_cases", "hdr_cases", "radius_cases"):
        meta[group] = man[group]
        for c in man[group]:
            for key in ("input_array", "output_array"):
                keep[c[key]] = z[c[key]]
    np.savez_compressed(os.path.join(HERE, "diffusion_cohort.npz"), **keep)
    with open(os.path.join(HERE, "diffusion_cohort.json"), "w") as fh:
        json.dump(meta, fh, indent=0)


BOUNDARY_TYPES = ("FjDiffusionSample", "FjDiffusionComponents", "FjDiffusionFailure", "FjDirBoundaryInput",
                  "FjDirBoundaryView", "FjDirBoundaryFailure")


def boundary_abi_facts() -> None:
    """Layouts of the diffusion-components and DIR-boundary import records,
    from the reference's C11 layout assertions (tests/ffi/prepared-values.c)
    -> boundary_abi_facts.json."""
    text = open(f"{REF}/tests/ffi/prepared-values.c").read()
    facts = {"types": {}, "fields": []}
    for name in BOUNDARY_TYPES:
        m = re.search(r"sizeof\(%s\) == (\d+) && _Alignof\(%s\) == (\d+)" % (name, name), text)
        facts["types"][name] = [int(m.group(1)), int(m.group(2))]
        for f, off in re.findall(r"offsetof\(%s, (\w+)\) == (\d+)" % name, text):
            facts["fields"].append([name, f, int(off)])
    json.dump(facts, open(os.path.join(HERE, "boundary_abi_facts.json"), "w"), indent=1)


if __name__ == "__main__":
    what = sys.argv[1:] or ["abi", "prepared", "lut", "dir", "halation", "diffusion"]
    if "abi" in what:
        abi_facts()
        abi_facts("prepared_call_abi_facts.inc", "call_abi_facts.json")
        boundary_abi_facts()
    if "prepared" in what:
        prepared_cases()
    if "lut" in what:
        scanner_luts()
    if "dir" in what:
        dir_fixtures()
    if "halation" in what:
        halation_fixtures()
    if "diffusion" in what:
        diffusion_fixtures()
